classdef FORWARD_SOLVER_CONVERGENT_BORN_MI355X < FORWARD_SOLVER_CONVERGENT_BORN
    % MI355X backend of the convergent Born series solver.
    % Same parameters and API as FORWARD_SOLVER_CONVERGENT_BORN_CUDA
    % (Codes/FORWARD_SOLVER_CONVERGENT_BORN_CUDA.m); the solve runs in
    % libcbs_hip.so through the CbsMi355x MEX gateway. All angles of a solve
    % call are one batched GPU solve.
    properties
       cbs_id;
    end
    methods(Static)
        function params=get_default_parameters(init_params)
            params=get_default_parameters@FORWARD_SOLVER_CONVERGENT_BORN();
            params.used_gpu=0;
            if nargin==1
                params=update_struct(params,init_params);
            end
        end
    end
    methods
        function h=FORWARD_SOLVER_CONVERGENT_BORN_MI355X(params)
            params.use_GPU=false; % the GPU part is done in the library
            h@FORWARD_SOLVER_CONVERGENT_BORN(params);
            g=struct;
            g.simulation_size=h.expected_RI_size(:)+2*h.boundary_thickness_pixel(:);
            g.RI_size=h.expected_RI_size(:);
            g.field_size=[params.size(1) params.size(2) 1]';
            if h.cyclic_boundary_xy
                g.conv_size=g.field_size;
            else
                g.conv_size=[g.RI_size(1:2)+g.field_size(1:2); 1];
            end
            g.used_gpu=uint32(params.used_gpu);
            g.RI_center=h.parameters.RI_center;
            g.pole_num=int32(h.pole_num);
            h.cbs_id=CbsMi355x(uint32(0),uint32(0),g,single(h.refocusing_util),single(h.kernel_trans),single(h.kernel_ref));
        end
        function set_RI(h,RI)
            if ~isequal(size(RI)',h.expected_RI_size(:))
                error(['The refractiv index does not have the expected size : ' ...
                    num2str(h.expected_RI_size(1)) ' ' num2str(h.expected_RI_size(2)) ' ' num2str(h.expected_RI_size(3))]);
            end
            set_RI@FORWARD_SOLVER_CONVERGENT_BORN(h,single(RI));
            p=struct;
            p.k0_nm=h.utility_border.k0_nm;
            p.eps_imag=h.eps_imag;
            p.fourier_res1=h.utility_border.fourier_space.res{1};
            p.fourier_res2=h.utility_border.fourier_space.res{2};
            p.fourier_res3=h.utility_border.fourier_space.res{3};
            p.ROI=h.ROI;
            p.Bornmax=h.Bornmax;
            p.dV=h.utility_border.dV;
            p.boundary_sharpness=h.parameters.boundary_sharpness;
            p.acyclic=double(h.parameters.acyclic);
            if h.cyclic_boundary_xy; p.shift_axes=4; else; p.shift_axes=7; end
            if h.boundary_thickness_pixel(1)==0 && h.boundary_thickness_pixel(2)==0; p.ramp_axes=4; else; p.ramp_axes=7; end
            CbsMi355x(uint32(1),uint32(h.cbs_id),p,single(h.V));
        end
        function delete(h)
            CbsMi355x(uint32(3),uint32(h.cbs_id));
        end
        function [fields_trans,fields_ref,fields_3D]=solve(h,input_field)
            input_field=single(input_field);
            if size(input_field,3)>1 && ~h.parameters.vector_simulation
                error('the source is 2D but parameter indicate a vectorial simulation');
            elseif size(input_field,3)==1 && h.parameters.vector_simulation
                error('the source is 3D but parameter indicate a non-vectorial simulation');
            end
            input_field=fftshift(fft2(ifftshift(input_field)));
            input_field=h.transform_field_3D(input_field);
            src=complex(fftshift(ifft2(ifftshift(input_field))));
            [t,r,u]=CbsMi355x(uint32(4),uint32(h.cbs_id),reshape(src,size(src,1),size(src,2),[]));
            A=size(input_field,4);
            P=size(input_field,3); % 1, or 3 field components for vector runs
            fields_trans=[]; fields_ref=[]; fields_3D=[];
            if h.parameters.return_3D
                fields_3D=reshape(u,size(u,1),size(u,2),size(u,3),P,A);
            end
            if h.parameters.return_transmission
                t=fftshift(fft2(ifftshift(reshape(t,size(t,1),size(t,2),P,A))));
                t=t+input_field;
                t=h.transform_field_2D(t);
                fields_trans=fftshift(ifft2(ifftshift(t)));
            end
            if h.parameters.return_reflection
                r=fftshift(fft2(ifftshift(reshape(r,size(r,1),size(r,2),P,A))));
                r=h.transform_field_2D_reflection(r);
                fields_ref=fftshift(ifft2(ifftshift(r)));
            end
        end
    end
end
